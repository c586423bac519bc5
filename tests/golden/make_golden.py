"""Writes the golden fixtures of tests/golden/ from data files of the reference's own tests.

  test_1p_cc_reference_p.txt  cell pressures of test/references/test_1p_cc-reference.vtu
                              (output of test_1p_incompressible_tpfa_numdiff: 10x10 unit square,
                              lens K 1e-12 in (0.2,0.8)^2, 1e-10 outside, Dirichlet p = 1e5 - 1e5 (y - 1)
                              on bottom/top, no-flow sides, gravity on; Float32 ascii)
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
    shutil.copy(os.path.join(REF, "test_pcsw_vangenuchten-reference.dat"), HERE)


if __name__ == "__main__":
    main()
