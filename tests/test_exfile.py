"""The golden-file readers and writer (iron_amd/exfile.py): exnode round trip,
derivatives and versions, an element-constant exelem field, and the nodal table
of an analytic-analysis report."""
from iron_amd import exfile


def test_exnode_round_trip(tmp_path):
    vals = {1: [[0.0], [0.5], [1.0], [-2.5e-3]], 2: [[20.0], [0.5], [1.0], [1.25e-17]]}
    p = tmp_path / "a.exnode"
    exfile.write_exnode(p, "Region", [("Geometry", "coordinate, rectangular cartesian", ["x", "y", "z"]),
                                      ("Pressure", "field, rectangular cartesian", ["1"])], vals)
    back = exfile.read_exnode(p)
    assert set(back) == {1, 2}
    assert back[2]["Geometry"] == [[20.0], [0.5], [1.0]] and back[2]["Pressure"] == [[1.25e-17]]


def test_exnode_derivatives_and_versions(tmp_path):
    text = """ Group name: Cube
 #Fields=1
 1) Dependent, field, rectangular cartesian, #Components=2
   1.  Value index= 1, #Derivatives= 1 (d/ds1), #Versions= 2
   2.  Value index= 5, #Derivatives= 0
 Node:            7
  1.0  2.0  3.0
  4.0
  5.0
"""
    p = tmp_path / "b.exnode"
    p.write_text(text)
    rec = exfile.read_exnode(p)[7]["Dependent"]
    assert rec == [[1.0, 2.0, 3.0, 4.0], [5.0]]


def test_exelem_element_constant_field(tmp_path):
    text = """ Group name: Cube
 Shape.  Dimension=3, line*line*line
 #Scale factor sets= 0
 #Nodes=           8
 #Fields=1
 1) Pressure, field, rectangular cartesian, #Components=1
  1.  l.Lagrange*l.Lagrange*l.Lagrange, no modify, grid based.
  #xi1=0, #xi2=0, #xi3=0
 Element:            1 0 0
   Values:
  -1.2500000000000000E+00
   Nodes:
            1           2           3           4           5           6           7           8
"""
    p = tmp_path / "c.exelem"
    p.write_text(text)
    el = exfile.read_exelem(p)
    assert el[1]["nodes"] == [1, 2, 3, 4, 5, 6, 7, 8]
    assert el[1]["grid"][("Pressure", 1)] == [-1.25]


def test_opanal_nodal_table(tmp_path):
    text = """ Analytic error analysis:

 Field 2 : Dependent Field

 Variable 1 : U

 Nodal errors:

   Node   Deriv     Numerical      Analytic        % error    Absolute error   Relative error
      1       1  0.10000E+01  0.10000E+01  0.0000E+00  0.0000E+00  0.0000E+00
      5       1  0.92113E+00  0.98231E+00  0.6228E+01  0.6118E-01  0.6228E-01

 RMS error =   0.4E-01

 Variable 2 : del U/del n

 Nodal errors:

      1       1 -0.50000E+00  0.00000E+00  0.0000E+00  0.5000E+00  0.5000E+00
"""
    p = tmp_path / "d.opanal"
    p.write_text(text)
    tab = exfile.read_opanal(p)
    assert tab["U"] == {1: [1.0, 1.0], 5: [0.92113, 0.98231]}
    assert tab["del U/del n"] == {1: [-0.5, 0.0]}
