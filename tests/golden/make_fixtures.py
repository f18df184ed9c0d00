"""Extract the reference's committed golden values into small JSON fixtures.

Run in a container that has the reference checkout:
    python tests/golden/make_fixtures.py /root/reference
Only numbers are extracted (nodal values of the reference's own output files);
no reference source is copied.
"""
import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def parse_opanal(path: Path):
    """Nodal blocks of an analytic-analysis output file: {variable: {node: (numerical, analytic)}}."""
    out, var = {}, None
    in_nodal = False
    for line in path.read_text().splitlines():
        m = re.match(r"\s*Variable\s+\d+\s*:\s*(.+)$", line)
        if m:
            var = m.group(1).strip()
            out[var] = {}
            in_nodal = False
            continue
        if line.strip().startswith("Nodal errors"):
            in_nodal = True
            continue
        if in_nodal:
            parts = line.split()
            if len(parts) >= 4 and parts[0].isdigit():
                out[var][int(parts[0])] = [float(parts[2]), float(parts[3])]
            elif line.strip().startswith("RMS"):
                in_nodal = False
    return out


def main(ref_root: str):
    ref = Path(ref_root)
    src = ref / "tests/ClassicalField/output_analytic_laplace/AnalyticLaplaceLinearLagrange.opanal"
    data = {
        "source": "tests/ClassicalField/output_analytic_laplace/AnalyticLaplaceLinearLagrange.opanal",
        "setup": {
            "comment": "tests/ClassicalField/AnalyticLaplace.F90:72-75,183-201,354-579: 2x2 bilinear Lagrange, "
                       "origin (-3.141592653579/2)^2, extent 2x2, analytic BCs u=cos(x)cosh(y) on boundary nodes",
            "n_elem": [2, 2], "origin_pi": 3.141592653579, "extent": [2.0, 2.0],
        },
        "nodal": parse_opanal(src),
    }
    (HERE / "analytic_laplace_linear_lagrange.json").write_text(json.dumps(data, indent=1))
    print("wrote analytic_laplace_linear_lagrange.json")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
