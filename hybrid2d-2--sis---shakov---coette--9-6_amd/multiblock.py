"""Meshes of several blocks and a driver that steps one GsisSolver handle per
block on one GPU.

The reference reads any number of blocks (generic.dat, src/sub_init.f90:102-178;
generic.inp, :236-321), links each inner subface to the matching subface of
its neighbour (Get_neighb_info, :636-693) and copies the neighbour's nodes into
the ghost nodes of inner faces (Update_coordinate_buffer_onemesh, :719-823).
Its kinetic step loops over the blocks (src/Hybrid_Sub.f90:29-40,180-187;
src/GKUA_DETFP.f90:33-70), and its kinetic BCs touch physical faces only
(src/GKUA_boundary.f90:25-35).  The ghost cells of an inner face are filled
from the neighbour only when a restart is read (GKUA_inner_mesco_boundary and
GKUA_inner_macro_boundary, src/Backup_write_read.f90:106-107); a fresh start
(init_flow_zero, src/sub_init.f90:459-462) never fills them.  Between two
restarts the blocks therefore step independently, each against the ghost
cells its inner faces last received.  MultiBlockSolver keeps that schedule:
resume() exchanges, initialize() and step() do not.

The RMS residual of a block is normalised by the cell count of the whole mesh
(src/GKUA_DETFP_RES.f90:54, Num_Cell summed over blocks at src/sub_init.f90:118)
and the monitor reports the maximum over blocks (src/IO.f90:231-235).
"""
from __future__ import annotations

import dataclasses
import pathlib
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import capi, restart
from .cases import BC, Case, ListDirected, bc_from_range, build_mesh, frd_shape, read_generic_dat_blocks  # noqa: F401


# ---------------------------------------------------------------------------
# the reference's input files with several blocks
# ---------------------------------------------------------------------------
def read_generic_inp_blocks(path: str | pathlib.Path, dims: Sequence[tuple]) -> List[List[BC]]:
    """generic.inp of read_bcin (src/sub_init.f90:236-321): a header record, NB,
    then per block NI NJ, the solver, the subface count and one record per
    subface (ist iend jst jend code).  A code of -1 is an inner face
    (negated to 1 on read, :278) followed by a record with the neighbour's
    range and block number (:282-284).  Only kinetic (GKUA, 2) blocks are
    supported.  ``dims`` holds (NI, NJ) of each block from generic.dat."""
    f = ListDirected(pathlib.Path(path).read_text())
    f.skip()
    nb = int(f.read(1)[0])
    if nb != len(dims):
        raise ValueError("block count in generic.inp differs from generic.dat")
    out = []
    for m in range(nb):
        ni, nj = (int(t) for t in f.read(2))
        if (ni, nj) != tuple(dims[m]):
            raise ValueError(f"block {m + 1}: generic.inp size differs from generic.dat")
        solver = int(f.read(1)[0])
        if solver != 2:
            raise NotImplementedError(f"block {m + 1}: only kinetic blocks (solver 2) are supported")
        nsub = int(f.read(1)[0])
        bcs = []
        for _ in range(nsub):
            ist, iend, jst, jend, code = (int(t) for t in f.read(5))
            b = bc_from_range(ist, iend, jst, jend, code, ni, nj)
            if b.code == 1:
                ist_n, iend_n, jst_n, jend_n, nbk = (int(t) for t in f.read(5))
                b.code = nbk
                b.nrange = (ist_n, iend_n, jst_n, jend_n)
            bcs.append(b)
        out.append(bcs)
    return out


def write_generic_dat_blocks(path: str | pathlib.Path, nodes: Sequence[tuple]) -> None:
    """generic.dat of several blocks (z = 0), readable by read_generic_dat_blocks
    and by the reference."""
    lines = [f"{len(nodes)}"]
    for xn, yn in nodes:
        ni, nj = xn.shape
        lines.append(f"{ni} {nj} 1")
        for a in (xn, yn, np.zeros_like(xn)):
            flat = np.asarray(a).ravel(order="F")
            for k in range(0, flat.size, 4):
                lines.append(" ".join(f"{v:.16e}" for v in flat[k:k + 4]))
    pathlib.Path(path).write_text("\n".join(lines) + "\n")


def write_generic_inp_blocks(path: str | pathlib.Path, dims: Sequence[tuple], bcs: Sequence[Sequence[BC]]) -> None:
    """generic.inp of several kinetic blocks; an inner face is written as code -1
    followed by its neighbour record."""
    lines = [" generic.inp", f"{len(dims)}"]
    for (ni, nj), bl in zip(dims, bcs):
        lines += [f"{ni} {nj}", "2", str(len(bl))]
        for b in bl:
            if b.uw != 0.0 or b.vw != 0.0:
                raise ValueError("generic.inp cannot express a moving wall")
            if b.code >= 1:
                if b.nrange is None:
                    raise ValueError("an inner face needs its neighbour range")
                lines.append(f"{b.ist:8d} {b.iend:8d} {b.jst:8d} {b.jend:8d} {-1:8d}")
                lines.append("".join(f"{v:8d} " for v in b.nrange) + f"{b.code:8d}")
            else:
                lines.append(f"{b.ist:8d} {b.iend:8d} {b.jst:8d} {b.jend:8d} {-b.code:8d}")
    pathlib.Path(path).write_text("\n".join(lines) + "\n")


# ---------------------------------------------------------------------------
# links between blocks
# ---------------------------------------------------------------------------
@dataclasses.dataclass(frozen=True)
class Link:
    """Inner subface ksub of block ``block`` reads subface msub of block
    ``nblock`` in orientation ``orient`` (2 same direction, 0 reversed).
    All numbers are 1-based."""
    block: int
    ksub: int
    nblock: int
    msub: int
    orient: int


def get_neighb_info(bcs: Sequence[Sequence[BC]]) -> List[Link]:
    """Get_neighb_info (src/sub_init.f90:636-693): for every inner subface, the
    neighbour subface whose range equals the neighbour range (orient 2) or
    equals it with i or j reversed (orient 0).  Every subface of the
    neighbour is tried and the last match is kept; no match is an error."""
    links = []
    for m, bl in enumerate(bcs, start=1):
        for ksub, b in enumerate(bl, start=1):
            if b.code < 1:
                continue
            if b.code > len(bcs):
                raise ValueError(f"block {m} subface {ksub}: neighbour block {b.code} does not exist")
            ib_n, ie_n, jb_n, je_n = b.nrange
            found = None
            for k2, c in enumerate(bcs[b.code - 1], start=1):
                if (c.ist, c.iend, c.jst, c.jend) == (ib_n, ie_n, jb_n, je_n):
                    found = (k2, 2)
                elif (c.ist, c.iend, c.jst, c.jend) == (ie_n, ib_n, jb_n, je_n):
                    found = (k2, 0)
                elif (c.ist, c.iend, c.jst, c.jend) == (ib_n, ie_n, je_n, jb_n):
                    found = (k2, 0)
            if found is None:
                raise ValueError("can not find the correct boundary for inner bound. condition "
                                 f"(block {m} subface {ksub})")
            links.append(Link(m, ksub, b.code, found[0], found[1]))
    return links


def build_block_meshes(nodes: Sequence[tuple], bcs: Sequence[Sequence[BC]], lref: float = 1.0):
    """Creat_Mesh + Update_coordinate_buffer_onemesh for every block: inner
    faces take their ghost nodes from the neighbour (src/sub_init.f90:739-770)."""
    links = get_neighb_info(bcs)
    meshes = []
    for m, ((xn, yn), bl) in enumerate(zip(nodes, bcs), start=1):
        inner = {}
        for ln in links:
            if ln.block == m:
                xnb, ynb = nodes[ln.nblock - 1]
                inner[ln.ksub - 1] = (xnb, ynb, bcs[ln.nblock - 1][ln.msub - 1], ln.orient)
        meshes.append(build_mesh(xn, yn, bl, lref, inner))
    return meshes, links


def split_nodes(xn: np.ndarray, yn: np.ndarray, bcs: Sequence[BC], m: int, axis: int = 0,
                rotate_second: bool = False):
    """Split one block at node index m (1-based) along i (axis 0) or j (axis 1)
    into two blocks that share that node line, with an inner face on each side
    of it; the physical BCs are clipped to each part.  With ``rotate_second``
    the second block's indices both run backwards (the block turned by 180
    degrees, still right-handed), so that the two inner faces are both high
    faces and meet in reversed orientation.  Returns (nodes, bcs) of the two
    blocks."""
    ni, nj = xn.shape
    n = (ni, nj)[axis]
    if not 3 <= m <= n - 2:
        raise ValueError("split node out of range")
    spans = ((1, m), (m, n))
    nodes, dims = [], []
    for k, (lo, hi) in enumerate(spans):
        sl = [slice(None), slice(None)]
        sl[axis] = slice(lo - 1, hi)
        a, b = xn[tuple(sl)], yn[tuple(sl)]
        if k == 1 and rotate_second:
            a, b = a[::-1, ::-1], b[::-1, ::-1]
        nodes.append((np.asfortranarray(a), np.asfortranarray(b)))
        dims.append(a.shape)

    def local(k, i, j):
        """node (i, j) of the whole block in block k's numbering"""
        if axis == 0:
            i = i - spans[k][0] + 1
        else:
            j = j - spans[k][0] + 1
        if k == 1 and rotate_second:
            i, j = dims[1][0] + 1 - i, dims[1][1] + 1 - j
        return i, j

    def as_range(k, p0, p1):
        (i0, j0), (i1, j1) = local(k, *p0), local(k, *p1)
        return (i0, i1, j0, j1)

    out = [[], []]
    for k, (lo, hi) in enumerate(spans):
        for b in bcs:
            r = [[b.ist, b.iend], [b.jst, b.jend]]
            a0, a1 = r[axis]
            if b.face in ((1, 3) if axis == 0 else (2, 4)):
                if not lo <= a0 <= hi:
                    continue
            else:
                a0, a1 = max(a0, lo), min(a1, hi)
                if a0 >= a1:
                    continue
            r[axis] = [a0, a1]
            i0, i1, j0, j1 = as_range(k, (r[0][0], r[1][0]), (r[0][1], r[1][1]))
            out[k].append(bc_from_range(min(i0, i1), max(i0, i1), min(j0, j1), max(j0, j1), -b.code,
                                        dims[k][0], dims[k][1], b.uw, b.vw))
    # the shared node line, from its first to its last node in the whole block
    p0, p1 = ((m, 1), (m, nj)) if axis == 0 else ((1, m), (ni, m))
    for k in (0, 1):
        i0, i1, j0, j1 = as_range(k, p0, p1)
        own = (min(i0, i1), max(i0, i1), min(j0, j1), max(j0, j1))
        flip = (i0, i1, j0, j1) != own  # this block runs along the line backwards
        o = 1 - k
        nr = as_range(o, p1, p0) if flip else as_range(o, p0, p1)
        b = bc_from_range(*own, -1, dims[k][0], dims[k][1])
        b.code, b.nrange = o + 1, nr
        out[k].append(b)
    return nodes, out


# ---------------------------------------------------------------------------
# driver
# ---------------------------------------------------------------------------
class MultiBlockSolver:
    """One solver handle per block.  ``make(case)`` builds a handle; by default
    a GsisSolver (the GPU path) with the keyword arguments given here."""

    def __init__(self, cases: Sequence[Case], links: Sequence[Link], make: Optional[Callable] = None,
                 mode: int = capi.MODE_M0, **kw):
        if mode != capi.MODE_M0 and len(cases) > 1:
            raise NotImplementedError("the synthetic NS solve of the reference works on block 1 only "
                                      "(src/Hybrid_Sub.f90:24,45-153): several blocks run in mode M0")
        if make is None:
            from .solver import GsisSolver

            def make(c):
                return GsisSolver(c, mode=mode, **kw)
        self.cases = list(cases)
        self.links = list(links)
        self.ncell = sum(c.ncell for c in self.cases)
        self.blocks = []
        for c in self.cases:
            self.blocks.append(make(c))

    def initialize(self):
        """init_flow_zero for every block; inner-face ghost cells keep the
        initial state, as in the reference."""
        for b in self.blocks:
            b.initialize()

    def exchange(self):
        """GKUA_inner_mesco_boundary + GKUA_inner_macro_boundary over all inner
        subfaces, blocks and subfaces in order."""
        for ln in self.links:
            self.blocks[ln.block - 1].inner_exchange(ln.ksub, self.blocks[ln.nblock - 1], ln.msub, ln.orient)

    def step(self, n: int = 1):
        for b in self.blocks:
            b.step(n)

    def residual(self) -> np.ndarray:
        """max over blocks of each block's RMS normalised by the whole mesh."""
        r = np.zeros(4)
        for c, b in zip(self.cases, self.blocks):
            r = np.maximum(r, b.residual() * np.sqrt(c.ncell / self.ncell))
        return r

    def counters(self):
        return self.blocks[0].counters()

    def frd(self) -> List[np.ndarray]:
        return [b.frd() for b in self.blocks]

    def run(self, max_steps: int, tol: float = 1e-8, show: int = 50):
        hist = []
        done = 0
        while done < max_steps:
            n = min(show, max_steps - done)
            self.step(n)
            done += n
            r = self.residual()
            hist.append((done, r.copy()))
            if np.all(r < tol):
                break
        return done, hist

    def write_output(self, directory, hybrid: bool = False):
        """The reference's Tecplot files, one zone per block."""
        from . import output
        return output.write_output(directory, [(c, b.macro()) for c, b in zip(self.cases, self.blocks)], hybrid)

    # -- restart (src/Backup_write_read.f90) -------------------------------
    def save_restart(self, root, flag: int):
        """backup_Macro_write(flag): one record per block, then flag.dat."""
        k, tt, _ = self.counters()
        return restart.write_restart(root, flag, self.frd(), tt, k)

    def resume(self, root) -> int:
        """init_flow_read + backup_Mesco_read (src/Backup_write_read.f90:56-109):
        every block's distribution, its kinetic BCs and moments, the saved step
        counter, then the inner-face exchange.  Returns the slot read."""
        frds, tt, kstep, slot = restart.read_restart(root, [frd_shape(c) for c in self.cases])
        for b, f in zip(self.blocks, frds):
            b.upload(frd=f)
            b.refresh_moments()
            b.set_counters(kstep, tt)
        self.exchange()
        return slot

    def close(self):
        for b in self.blocks:
            b.close()


def block_cases(nodes: Sequence[tuple], bcs: Sequence[Sequence[BC]], template: Case, lref: float = 1.0):
    """Cases of several blocks that share ``template``'s velocity grid and constants."""
    meshes, links = build_block_meshes(nodes, bcs, lref)
    cases = [Case(f"{template.name}_b{m}", mesh, list(bl), template.vi, template.wi, template.vj, template.wj,
                  dict(template.const), template.init)
             for m, (mesh, bl) in enumerate(zip(meshes, bcs), start=1)]
    return cases, links
