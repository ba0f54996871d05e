"""Writes the golden fixtures of the GSIS tests from the reference's data files.

Run once in the build container (the reference tree is not present on GPU
machines):  python tests/golden/make_fixtures.py /root/reference/Hybrid2D

Fixtures are data only: the reference's input files for its two shipped
cases, and known-answer values read from its recorded output files and from
the numeric tables in its sources.
"""
import hashlib
import json
import pathlib
import re
import shutil
import sys

HERE = pathlib.Path(__file__).resolve().parent


def main(ref: pathlib.Path):
    inp = ref / "input"
    # 1) the two shipped cases' inputs (unchanged formats)
    cyl = HERE / "cylinder"
    cyl.mkdir(exist_ok=True)
    for f in ("generic.dat", "generic.inp", "control.in", "Quadrature_Para.dat"):
        shutil.copyfile(inp / f, cyl / f)
    ch = HERE / "channel"
    ch.mkdir(exist_ok=True)
    sub = [p for p in inp.iterdir() if p.is_dir()][0]  # the 50x6 channel case folder
    for f in ("generic.dat", "generic.inp"):
        shutil.copyfile(sub / f, ch / f)
    for f in ("control.in", "Quadrature_Para.dat"):
        shutil.copyfile(inp / f, ch / f)

    # 2) output/flow2d-GKUA.dat: moments of the uniform Ma=3 Maxwellian on the
    #    60x60 GL grid, written by output() (src/IO.f90:370-389), 41x41 cells
    lines = (ref / "output" / "flow2d-GKUA.dat").read_text().splitlines()
    rows = [list(map(float, ln.split())) for ln in lines[2:] if ln.strip()]
    cols = "x,y,d,u,v,T,p,txx,tyy,txy,qx,qy,tzz,cf".split(",")
    kat = {"columns": cols, "ni": 41, "nj": 41,
           "x1": [r[0] for r in rows], "y1": [r[1] for r in rows],
           "moments": {c: rows[0][k] for k, c in enumerate(cols) if k >= 2},
           "all_rows_equal": all(r[2:] == rows[0][2:] for r in rows)}
    (HERE / "kat_flow2d_gkua.json").write_text(json.dumps(kat))

    # 3) output/flow2d.dat first zone: initial NS fields on cells 3..n-3
    lines = (ref / "output" / "flow2d.dat").read_text().splitlines()
    m = re.search(r"i=\s*(\d+)\s*j=\s*(\d+)", lines[1])
    ni, nj = int(m.group(1)), int(m.group(2))
    rows = [list(map(float, ln.split())) for ln in lines[2:2 + ni * nj]]
    (HERE / "kat_flow2d.json").write_text(json.dumps(
        {"ni": ni, "nj": nj, "columns": "x,y,d,u,v,T,p,Ma".split(","), "rows": rows}))

    # 4) the GAULEG table (src/GKUA_lisansudu.f90:298-353) as decimal strings
    src = (ref / "src" / "GKUA_lisansudu.f90").read_text(errors="replace").splitlines()
    table, kgln = {}, None
    for ln in src[297:353]:
        mk = re.search(r"KGLN\.EQ\.(\d+)", ln)
        if mk:
            kgln = mk.group(1)
            table[kgln] = {"AGL": {}, "HGL": {}}
        for name, k, val in re.findall(r"(AGL|HGL)\((\d+)\)=([0-9.Ee+-]+)", ln):
            table[kgln][name][k] = val
    (HERE / "gauleg_table.json").write_text(json.dumps(table, indent=1))

    # 4b) the half-range Gauss-Hermite table GaussHermite
    #     (src/GKUA_lisansudu.f90:364-436) as decimal strings
    table, nv = {}, None
    for ln in src[363:436]:
        mk = re.search(r"NV\.EQ\.(\d+)", ln)
        if mk:
            nv = mk.group(1)
            table[nv] = {"AGH": {}, "HGH": {}}
        for name, k, val in re.findall(r"(AGH|HGH)\((\d+)\)=([0-9.Ee+-]+)", ln):
            table[nv][name][k] = val
    (HERE / "gauher_table.json").write_text(json.dumps(table, indent=1))

    # 5) fort.21: the constants line printed by init (src/sub_init.f90:62-64)
    txt = (ref / "fort.21").read_text().split()
    nums = [float(t) for t in txt if re.match(r"^-?[0-9.]+(E[+-]?\d+)?$", t)]
    (HERE / "fort21.json").write_text(json.dumps({"KHS": nums[0], "Ma": nums[1], "Re": nums[2],
                                                  "Kn_inf": nums[3], "Rcon": nums[4]}))

    # 6) SHA-256 and size of the output files that the shipped cylinder's
    #    initial state determines (output, src/IO.f90:307-553); the tests
    #    write them from the oracle's and the GPU's fields and compare bytes
    names = ("flow2d.dat", "flow2d-GKUA.dat", "flow2d-S0.dat", "Tao_g_i.dat", "Tao_g_j.dat", "q_g_i.dat",
             "q_g_j.dat", "flow2d_Dem.dat")
    dig = {}
    for n in names:
        raw = (ref / "output" / n).read_bytes()
        dig[n] = {"sha256": hashlib.sha256(raw).hexdigest(), "bytes": len(raw), "lines": raw.count(b"\n")}
    (HERE / "output_digests.json").write_text(json.dumps(dig, indent=1))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main(pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/Hybrid2D"))
