"""Writes the golden fixtures of tests/golden/ from data files of the reference's own tests.

  test_1p_cc_reference_p.txt  cell pressures of test/references/test_1p_cc-reference.vtu
                              (output of test_1p_incompressible_tpfa_numdiff: 10x10 unit square,
                              lens K 1e-12 in (0.2,0.8)^2, 1e-10 outside, Dirichlet p = 1e5 - 1e5 (y - 1)
                              on bottom/top, no-flow sides, gravity on; Float32 ascii)
  test_2p_incompressible_cc_reference.txt
                              cell data of test/references/test_2p_incompressible_cc-reference.vtu
                              (7th output of test_2p_incompressible_tpfa / _tpfa_analytic: 48x32 cells
                              on (0,6)x(0,4), DNAPL lens, adaptive dt from 250 s; Float32 ascii):
                              columns S_aq p_aq S_napl p_napl pc mob_aq mob_napl, cells x-fastest
  test_pcsw_vangenuchten-reference.dat is copied verbatim from test/references/.

Run from the repository root with the reference tree at /root/reference. Only data is written.
"""
import os
import re
import shutil

REF = "/root/reference/test/references"
HERE = os.path.dirname(os.path.abspath(__file__))


def cell_array(vtu_text, name):
    m = re.search(r'<DataArray[^>]*Name="%s"[^>]*format="ascii"[^>]*>(.*?)</DataArray>' % re.escape(name),
                  vtu_text, re.S)
    return [float(v) for v in m.group(1).split()]


def main():
    with open(os.path.join(REF, "test_1p_cc-reference.vtu")) as f:
        p = cell_array(f.read(), "p")
    assert len(p) == 100
    with open(os.path.join(HERE, "test_1p_cc_reference_p.txt"), "w") as f:
        f.write("# cell pressures [Pa] of test/references/test_1p_cc-reference.vtu (cells x-fastest)\n")
        for v in p:
            f.write(f"{v:.9g}\n")
    names = ["S_aq", "p_aq", "S_napl", "p_napl", "pc", "mob_aq", "mob_napl"]
    with open(os.path.join(REF, "test_2p_incompressible_cc-reference.vtu")) as f:
        text = f.read()
    cols = [cell_array(text, nm) for nm in names]
    assert all(len(c) == 48 * 32 for c in cols)
    with open(os.path.join(HERE, "test_2p_incompressible_cc_reference.txt"), "w") as f:
        f.write("# cell data of test/references/test_2p_incompressible_cc-reference.vtu (cells x-fastest): "
                + " ".join(names) + "\n")
        for row in zip(*cols):
            f.write(" ".join(f"{v:.9g}" for v in row) + "\n")
    shutil.copy(os.path.join(REF, "test_pcsw_vangenuchten-reference.dat"), HERE)


if __name__ == "__main__":
    main()
