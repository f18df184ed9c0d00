"""Extract the reference's committed golden values into small JSON fixtures.

Run in a container that has the reference checkout:
    python tests/golden/make_fixtures.py /root/reference
Only numbers are extracted (nodal values of the reference's own output files);
no reference source is copied.
"""
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
from iron_amd.exfile import read_exelem, read_exnode, read_opanal  # noqa: E402


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
        "nodal": read_opanal(src),
    }
    (HERE / "analytic_laplace_linear_lagrange.json").write_text(json.dumps(data, indent=1))
    print("wrote analytic_laplace_linear_lagrange.json")

    # Finite-elasticity cantilevers (Lagrange, element-based pressure): nodal deformed
    # coordinates and DELUDELN values, element pressures and their DELUDELN values.
    cases = {
        "mr_cantilever_python_1x1x1.json": (
            "bindings/python/tests/output_cantilever/Cantilever.part0",
            {"comment": "bindings/python/tests/Cantilever.py: 1x1x1 trilinear, extent 60x40x40, "
                        "element pressure (initial -8), c1=2, c2=2, density 9e-4, gravity (0,0,-9.81), "
                        "2 load increments, nodes 1,3,5,7 fixed (AddNode FIXED 0)",
             "n_elem": [1, 1, 1], "extent": [60.0, 40.0, 40.0], "c1": 2.0, "c2": 2.0,
             "density": 9e-4, "gravity": [0.0, 0.0, -9.81], "increments": 2, "pressure0": -8.0}),
        "mr_cantilever_linear_3x2x2.json": (
            "tests/FiniteElasticity/output_cantilever/linear_3x2x2/Cantilever.part0",
            {"comment": "tests/FiniteElasticity/Cantilever.F90 (defaults, :150-153, :185-206): 3x2x2 "
                        "trilinear, extent 60x40x40, element pressure (initial -14), c1=2, c2=6, "
                        "density 9e-4, gravity (0,0,-9.8), 2 load increments, x=0 face fixed",
             "n_elem": [3, 2, 2], "extent": [60.0, 40.0, 40.0], "c1": 2.0, "c2": 6.0,
             "density": 9e-4, "gravity": [0.0, 0.0, -9.8], "increments": 2, "pressure0": -14.0}),
    }
    for name, (stem, setup) in cases.items():
        nodes = read_exnode(ref / (stem + ".exnode"))
        elems = read_exelem(ref / (stem + ".exelem"))
        data = {
            "source": stem + ".exnode/.exelem",
            "setup": setup,
            "nodes": {n: {"dependent": [rec["Dependent"][c][0] for c in range(3)],
                          "deludeln": [rec["del U/del n"][c][0] for c in range(3)]}
                      for n, rec in nodes.items()},
            "elements": {e: {"nodes": rec["nodes"],
                             "pressure": rec["grid"][("Dependent", 4)][0],
                             "deludeln_p": rec["grid"][("del U/del n", 4)][0]}
                         for e, rec in elems.items()},
        }
        (HERE / name).write_text(json.dumps(data, indent=1))
        print("wrote", name)

    # Nearly incompressible Mooney-Rivlin simple shear (no pressure field): nodal
    # deformed coordinates and DELUDELN values, element connectivity.
    stem = "tests/FiniteElasticity/output_shear/SimpleShear.part0"
    nodes = read_exnode(ref / (stem + ".exnode"))
    elems = read_exelem(ref / (stem + ".exelem"))
    data = {
        "source": stem + ".exnode/.exelem",
        "setup": {"comment": "tests/FiniteElasticity/SimpleShear.F90: 2x2x2 triquadratic Lagrange, 3x3x3 Gauss, "
                             "extent 1x1x1, EQUATIONS_SET_NEARLY_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE (3 "
                             "components), materials c1=0.5, c2=0, kappa=1e4, top (z=1) nodes fixed at "
                             "(x+0.1, y, z), bottom (z=0) nodes fixed at (x, y, z), one load step, Newton with "
                             "the equations-calculated (finite-difference) Jacobian and a direct linear solve",
                  "n_elem": [2, 2, 2], "extent": [1.0, 1.0, 1.0], "c1": 0.5, "c2": 0.0, "kappa": 10000.0,
                  "shear": 0.1},
        "nodes": {n: {"dependent": [rec["Dependent"][c][0] for c in range(3)],
                      "deludeln": [rec["del U/del n"][c][0] for c in range(3)]}
                  for n, rec in nodes.items()},
        "elements": {e: {"nodes": rec["nodes"]} for e, rec in elems.items()},
    }
    (HERE / "mr_simple_shear_2x2x2.json").write_text(json.dumps(data, indent=1))
    print("wrote mr_simple_shear_2x2x2.json")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
