"""Write libxc's own LDA_X / LDA_C_PZ regression data as a JSON fixture.

Reads only data files of the vendored libxc testsuite (no code is run):
  ExLib/libxc/testsuite/input/{H,Li,BrOH,BrOH+}           (rho_a rho_b ... per row)
  ExLib/libxc/testsuite/regression/lda_x/lda_x.<sys>.unpol.{0,1}.bz2
  ExLib/libxc/testsuite/regression/lda_c/lda_c_pz.<sys>.unpol.{0,1}.bz2
  ExLib/libxc/testsuite/df_repo/{lda_x,lda_c_pz}.data     (rho_a == rho_b rows)
Unpolarised inputs are rho = rho_a + rho_b (testsuite/xc-regression.c:249-261).
Order-0 files hold eps (energy per particle); order-1 files hold v = d(n eps)/dn.
df_repo ``zk`` is the energy per volume n*eps (12 significant digits).
"""
import bz2
import json
import os
import re
import sys

REF = "/root/reference/ExLib/libxc/testsuite"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libxc_lda.json")


def read_input(system):
    with open(os.path.join(REF, "input", system)) as f:
        n = int(f.readline())
        rows = [list(map(float, f.readline().split())) for _ in range(n)]
    return [r[0] + r[1] for r in rows]


def read_reg(path):
    with bz2.open(path, "rt") as f:
        head = f.readline().split()
        f.readline()
        vals = [float(f.readline()) for _ in range(int(head[1]))]
    return vals


def read_df_repo(name):
    out = []
    txt = open(os.path.join(REF, "df_repo", name + ".data")).read()
    for block in txt.split("rhoa=")[1:]:
        ra = float(block.split()[0])
        rb = float(re.search(r"rhob=\s*(\S+)", block).group(1))
        zk = float(re.search(r"zk\s*=\s*(\S+)", block).group(1))
        va = float(re.search(r"vrhoa\s*=\s*(\S+)", block).group(1))
        if ra == rb:
            out.append({"rho": ra + rb, "zk": zk, "vrho": va})
    return out


def main():
    fix = {"regression": {}, "df_repo": {}}
    for func, sub in (("lda_x", "lda_x"), ("lda_c_pz", "lda_c")):
        for system in ("H", "Li", "BrOH", "BrOH+"):
            rho = read_input(system)
            base = os.path.join(REF, "regression", sub, f"{func}.{system}.unpol")
            fix["regression"][f"{func}.{system}"] = {
                "rho": rho, "exc": read_reg(base + ".0.bz2"), "vxc": read_reg(base + ".1.bz2")}
        fix["df_repo"][func] = read_df_repo(func)
    with open(OUT, "w") as f:
        json.dump(fix, f, indent=0)
    print("wrote", OUT, file=sys.stderr)


if __name__ == "__main__":
    main()
