"""Several blocks on one GPU, through the C-ABI.

gsis_inner_exchange is checked bit for bit against the oracle's restatement
of GKUA_inner_mesco_boundary + GKUA_inner_macro_boundary (itself pinned by the
split-block geometry tests on the CPU).  Two-block cavities, a cavity with a
symmetry face and a multi-block restart are checked against oracle blocks
driven the same way, to the FP64 bar of 1e-10 relative L2."""
import numpy as np
import pytest

from helpers import load_oracle, load_pkg, rel_l2, solver_mod

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _split_cavity(n=16, nv=16, kn=0.1, axis=0, rot=False, lid=0.15, dt=0.5, scheme=1):
    g = load_pkg()
    whole = g.cases.cavity_case(n, nv, kn=kn, lid=lid, dt=dt)
    whole.const["scheme"] = scheme
    nx, ny = whole.nx, whole.ny
    xn = np.asfortranarray(whole.mesh.x[2:nx + 2, 2:ny + 2])
    yn = np.asfortranarray(whole.mesh.y[2:nx + 2, 2:ny + 2])
    m = (nx if axis == 0 else ny) // 2 + 1
    nodes, bcs = g.multiblock.split_nodes(xn, yn, whole.bcs, m, axis, rot)
    return g.multiblock.block_cases(nodes, bcs, whole)


def _oracle_blocks(cases, links):
    orc = load_oracle()
    return load_pkg().multiblock.MultiBlockSolver(cases, links, make=lambda c: orc.OracleSolver(c))


def _compare_blocks(gpu, cpu, tag, residual=True):
    for k, (a, b) in enumerate(zip(gpu.blocks, cpu.blocks)):
        ma, mc = a.macro(), b.macro()
        errs = {"frd": rel_l2(a.frd(), b.frd()), "s": rel_l2(ma["s"], mc["s"]), "cf": rel_l2(ma["cf"], mc["cf"]),
                "swall": rel_l2(ma["swall"], mc["swall"])}
        print(tag, "block", k + 1, {q: f"{v:.1e}" for q, v in errs.items()}, flush=True)
        bad = {q: v for q, v in errs.items() if not v <= TOL}
        assert not bad, f"{tag} block {k + 1}: {bad}"
    if residual:  # kept by a step: none before the first
        rg, rc = gpu.residual(), cpu.residual()
        assert rel_l2(rg, rc) <= TOL, (rg, rc)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
@pytest.mark.parametrize("axis,rot", [(0, False), (0, True), (1, False), (1, True)])
def test_inner_exchange_bitwise(prec, axis, rot):
    """FRD(1:2) of every velocity and S_GKUA(1:10) of two cell layers, in the
    orientation Get_neighb_info found; nothing else changes."""
    g = load_pkg()
    orc = load_oracle()
    cases, links = _split_cavity(n=12, nv=8, axis=axis, rot=rot)
    print("exchange", "fp32" if prec else "fp64", "axis", axis, "rot", rot, links, flush=True)
    mbs = g.multiblock.MultiBlockSolver(cases, links, precision=prec)
    rng = np.random.default_rng(7 + axis + 2 * rot)
    for c, b in zip(cases, mbs.blocks):
        frd = np.asfortranarray(rng.random(g.cases.frd_shape(c)) + 0.5)
        s = np.asfortranarray(rng.random((11, c.nx + 7, c.ny + 7)) + 0.5)
        b.upload(frd=frd, s=s)
    pre = [(b.frd(), b.macro()["s"]) for b in mbs.blocks]
    exp = [(f.copy(), s.copy()) for f, s in pre]
    for ln in links:
        fd, sd = exp[ln.block - 1]
        fs, ss = pre[ln.nblock - 1]
        orc.inner_exchange_arrays(fd, sd, cases[ln.block - 1].bcs[ln.ksub - 1], fs, ss,
                                  cases[ln.nblock - 1].bcs[ln.msub - 1], ln.orient)
    mbs.exchange()
    for k, b in enumerate(mbs.blocks):
        np.testing.assert_array_equal(b.frd(), exp[k][0])
        np.testing.assert_array_equal(b.macro()["s"], exp[k][1])
        assert not np.array_equal(exp[k][0], pre[k][0])
    mbs.close()


@pytest.mark.parametrize("axis,rot,scheme", [(0, False, 1), (1, True, 1), (0, True, 3), (1, False, 0)],
                         ids=["i-nnd", "j-turned-nnd", "i-turned-weno3", "j-up1"])
def test_two_block_cavity_parity(axis, rot, scheme):
    """Both blocks step independently between exchanges, as between two
    restarts of the reference; the GPU blocks follow the oracle blocks, with
    each flux scheme."""
    g = load_pkg()
    cases, links = _split_cavity(n=16, nv=16, axis=axis, rot=rot, scheme=scheme)
    print("two-block cavity axis", axis, "rot", rot, "scheme", scheme, [c.name for c in cases], flush=True)
    gpu = g.multiblock.MultiBlockSolver(cases, links)
    cpu = _oracle_blocks(cases, links)
    for s in (gpu, cpu):
        s.initialize()
    _compare_blocks(gpu, cpu, "init", residual=False)
    for rnd in range(2):
        for s in (gpu, cpu):
            s.exchange()
            s.step(8)
        _compare_blocks(gpu, cpu, f"round {rnd + 1}")
    gpu.close()
    cpu.close()


def test_symmetry_face_parity():
    """A symmetry face is passive in the kinetic step: its ghost cells keep the
    initial state, and the fluxes through it use them unchanged."""
    g = load_pkg()
    cs, capi = g.cases, g.capi
    base = cs.cavity_case(12, 16, kn=0.1, lid=0.15, dt=0.5)
    bcs = list(base.bcs)
    b = bcs[1]  # i = nx
    bcs[1] = cs.BC(b.face, b.ist, b.iend, b.jst, b.jend, capi.BC_SYMMETRY)
    case = cs.Case("cavity12_sym", base.mesh, bcs, base.vi, base.wi, base.vj, base.wj, base.const, "rest")
    print("case:", case.name, flush=True)
    gpu = solver_mod().GsisSolver(case)
    cpu = load_oracle().OracleSolver(case)
    for s in (gpu, cpu):
        s.initialize()
    f0 = gpu.frd()
    gpu.step(15)
    cpu.step(15)
    fg = gpu.frd()
    assert rel_l2(fg, cpu.frd()) <= TOL
    assert rel_l2(gpu.macro()["s"], cpu.macro()["s"]) <= TOL
    assert rel_l2(gpu.residual(), cpu.residual()) <= TOL
    nx = case.nx
    np.testing.assert_array_equal(fg[:, nx + 3:nx + 5, 4:-4], f0[:, nx + 3:nx + 5, 4:-4])
    gpu.close()
    cpu.close()


def test_multiblock_backup_and_resume(tmp_path):
    """save_restart writes one record per block; resume() reads them, refreshes
    BCs and moments, restores the counters and exchanges the inner faces
    (backup_Mesco_read, src/Backup_write_read.f90:93-109); the resumed blocks
    then follow oracle blocks resumed the same way."""
    g = load_pkg()
    orc = load_oracle()
    cases, links = _split_cavity(n=14, nv=12, axis=0, rot=True)
    print("multi-block restart", [c.name for c in cases], flush=True)
    a = g.multiblock.MultiBlockSolver(cases, links)
    a.initialize()
    a.step(6)
    a.save_restart(tmp_path, 1)
    saved = a.frd()
    a.close()
    b = g.multiblock.MultiBlockSolver(cases, links)
    b.initialize()
    assert b.resume(tmp_path) == 1
    assert b.counters()[0] == 6
    exp = [f.copy() for f in saved]
    for ln in links:
        c1, c2 = cases[ln.block - 1], cases[ln.nblock - 1]
        orc.inner_exchange_arrays(exp[ln.block - 1], np.zeros((11, c1.nx + 7, c1.ny + 7)), c1.bcs[ln.ksub - 1],
                                  saved[ln.nblock - 1], np.zeros((11, c2.nx + 7, c2.ny + 7)), c2.bcs[ln.msub - 1],
                                  ln.orient)
    for f, e in zip(b.frd(), exp):
        # cells -2..n+2: all but the wall slot, which the refreshed BCs rewrite
        np.testing.assert_array_equal(f[:, 1:-1, 1:-1], e[:, 1:-1, 1:-1])
    cpu = _oracle_blocks(cases, links)
    cpu.initialize()
    for blk, f in zip(cpu.blocks, saved):
        blk.upload(frd=f)
        blk.refresh_moments()
    cpu.exchange()
    b.step(5)
    cpu.step(5)
    _compare_blocks(b, cpu, "resumed +5")
    b.close()
    cpu.close()


def test_inner_exchange_argument_errors():
    """gsis_inner_exchange refuses, with GSIS_EARG and before any device work,
    a receiving subface that is not an inner face, subface numbers out of
    range, faces of different cell counts and blocks of different precisions."""
    g = load_pkg()
    capi = g.capi
    cases, links = _split_cavity(n=12, nv=8, axis=0, rot=False)
    ln = links[0]
    a = g.multiblock.MultiBlockSolver(cases, links)
    b32 = solver_mod().GsisSolver(cases[1], precision=1)
    dst, src = a.blocks[ln.block - 1], a.blocks[ln.nblock - 1]
    wall = [k for k, c in enumerate(cases[ln.block - 1].bcs, start=1) if c.code < 1][0]
    other_face = [k for k, c in enumerate(cases[ln.nblock - 1].bcs, start=1)
                  if (c.iend - c.ist + c.jend - c.jst) != (cases[ln.block - 1].bcs[ln.ksub - 1].iend
                                                          - cases[ln.block - 1].bcs[ln.ksub - 1].ist
                                                          + cases[ln.block - 1].bcs[ln.ksub - 1].jend
                                                          - cases[ln.block - 1].bcs[ln.ksub - 1].jst)]
    bad = [(dst, wall, src, ln.msub), (dst, 0, src, ln.msub), (dst, ln.ksub, src, 99), (dst, ln.ksub, b32, ln.msub)]
    if other_face:
        bad.append((dst, ln.ksub, src, other_face[0]))
    for d, k, s_, m in bad:
        with pytest.raises(capi.GsisError) as e:
            d.inner_exchange(k, s_, m, ln.orient)
        assert e.value.code == capi.GSIS_EARG, str(e.value)
    a.exchange()  # the handles are still usable
    b32.close()
    a.close()
