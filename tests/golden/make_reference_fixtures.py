"""Extract the reference's golden arrays (test DATA, not code) into reference_fixtures.json.

Run in the build container only (it reads the reference checkout); the JSON it writes is
committed and is all that tests read at run time:

    python tests/golden/make_reference_fixtures.py /root/reference

Sources (relative to the reference root):
  tests/basic/MeshQuad4.cpp:10-100      Regular(5,3): coor, conn, dofs, node sets, periodic
  tests/basic/MeshQuad4.cpp:103-256     FineLayer(9,18): coor, conn, dofs, node sets, periodic
  tests/basic/MeshTri3.cpp:6-88         Tri3 Regular(5,3): coor, conn, dofs, node sets, periodic
  tests/basic/MeshHex8.cpp:8-188        Regular(3,3,3): coor, conn, node sets, dofs, periodic
  tests/basic/MeshHex8.cpp:192-377      FineLayer(3,3,3): coor, conn, node sets, dofs, periodic
  tests/basic/Mesh.cpp:199-236          Regular(3,3): coordination, elem2node
  tests/basic/Mesh.cpp:207-213, 236-258 Regular(2,2,h=2) centers; Regular(3,3) elemmap2nodemap
  tests/basic/Mesh.cpp:260-382          FineLayer(3,3) roll + elemmap2nodemap, examples 1 and 2
  tests/test_Mesh.py:23-171             Regular(4,4) periodic: coordination, node2dof,
                                        nodaltyings, elem2node(conn, dofsPeriodic)
  tests/basic/MatrixDiagonal.cpp:12-53  lumped mass of Regular(2,2) (Nodal quadrature)
  tests/basic/Vector.cpp:11-167         periodic AsDofs inputs and expected index pairs
  tests/basic/VectorPartitioned.cpp:10-38          Regular(3,3), iip = arange(8): dofs_is_p
  tests/basic/MatrixDiagonalPartitioned.cpp:12-53  Regular(2,2), iip = {0, 2}: lumped mass data()
"""

from __future__ import annotations

import ast
import json
import os
import re
import sys


def _braced(src: str, start: int) -> str:
    """Return the {...} literal that starts at src[start] == '{'."""
    depth = 0
    for i in range(start, len(src)):
        if src[i] == "{":
            depth += 1
        elif src[i] == "}":
            depth -= 1
            if depth == 0:
                return src[start : i + 1]
    raise ValueError("unbalanced braces")


def _cpp_literal(text: str):
    body = text.replace("{", "[").replace("}", "]")
    body = re.sub(r"(\d)\.(?=[\s,\]])", r"\1.0", body)
    return ast.literal_eval(body)


def cpp_arrays(src: str) -> dict:
    """All `<type> name_ = {...};` and `size_t name_ = N;` definitions in a section."""
    out = {}
    for m in re.finditer(r"xt::xtensor<[^>]+>\s+(\w+)\s*=\s*\{", src):
        out[m.group(1)] = _cpp_literal(_braced(src, m.end() - 1))
    for m in re.finditer(r"size_t\s+(\w+_)\s*=\s*(\d+)\s*;", src):
        out[m.group(1)] = int(m.group(2))
    return out


def section(src: str, name: str) -> str:
    i = src.index('SECTION("%s")' % name)
    j = src.find("SECTION(", i + 10)
    return src[i : j if j > 0 else len(src)]


def py_function_literal(path: str, func: str):
    """The first list literal assigned to `real` inside `func` of a unittest file."""
    tree = ast.parse(open(path).read())
    for node in ast.walk(tree):
        if isinstance(node, ast.FunctionDef) and node.name == func:
            for st in ast.walk(node):
                if isinstance(st, ast.Assign) and getattr(st.targets[0], "id", None) == "real":
                    v = st.value
                    if isinstance(v, ast.Call):  # np.array([...]) possibly .ravel()
                        while isinstance(v, ast.Call) and not (
                            isinstance(v.func, ast.Attribute) and v.func.attr == "array"
                        ):
                            v = v.func.value
                        v = v.args[0]
                    return ast.literal_eval(v)
    raise KeyError(func)


def main(root: str) -> None:
    basic = os.path.join(root, "tests", "basic")
    fx: dict = {}

    src = open(os.path.join(basic, "MeshQuad4.cpp")).read()
    fx["MeshQuad4_Regular_5_3"] = cpp_arrays(section(src, "Regular"))

    fx["MeshQuad4_FineLayer_9_18"] = cpp_arrays(section(src, "FineLayer"))

    src = open(os.path.join(basic, "MeshTri3.cpp")).read()
    fx["MeshTri3_Regular_5_3"] = cpp_arrays(section(src, "Regular"))

    src = open(os.path.join(basic, "MeshHex8.cpp")).read()
    fx["MeshHex8_Regular_3_3_3"] = cpp_arrays(section(src, "Regular"))
    fx["MeshHex8_FineLayer_3_3_3"] = cpp_arrays(section(src, "FineLayer - basic"))

    src = open(os.path.join(basic, "Mesh.cpp")).read()
    sec = section(src, "coordination")
    fx["Mesh_coordination_Regular_3_3"] = cpp_arrays(sec)["ret"]
    fx["Mesh_centers_Regular_2_2_h2"] = cpp_arrays(section(src, "centers"))["c"]
    fx["Mesh_elemmap2nodemap_Regular_3_3"] = cpp_arrays(section(src, "elemmap2nodemap"))
    for k in (1, 2):
        fx[f"Mesh_elemmap2nodemap_FineLayer_3_3_example{k}"] = cpp_arrays(
            section(src, f"elemmap2nodemap - example {k}"))
    sec = section(src, "elem2node")
    lists = re.findall(r"tonode\[(\d+)\] == std::vector<size_t>\{([^}]*)\}", sec)
    fx["Mesh_elem2node_Regular_3_3"] = [
        [int(v) for v in body.split(",")] for _, body in sorted(lists, key=lambda t: int(t[0]))
    ]

    tm = os.path.join(root, "tests", "test_Mesh.py")
    fx["test_Mesh_coordination_dofs_4_4"] = py_function_literal(tm, "test_coordination_dofs")
    fx["test_Mesh_node2dof_4_4"] = py_function_literal(tm, "test_coordination_node2dof")
    fx["test_Mesh_nodaltyings_4_4"] = py_function_literal(tm, "test_coordination_nodaltyings")
    fx["test_Mesh_elem2node_periodic_4_4"] = py_function_literal(tm, "test_elem2node_periodic")

    src = open(os.path.join(basic, "MatrixDiagonal.cpp")).read()
    sec = section(src, "assemble")
    m = re.search(r"xt::xtensor<double, 1> a = \{", sec)
    fx["MatrixDiagonal_lumped_mass_Regular_2_2"] = _cpp_literal(_braced(sec, m.end() - 1))

    src = open(os.path.join(basic, "Vector.cpp")).read()
    vec = {}
    for name in ("asDofs - nodevec", "asDofs - assembleDofs"):
        sec = section(src, name)
        assigns = re.findall(r"\b([vf])\((\d+), (\d+)\) = ([-\d.]+);", sec)
        arr = [[0.0, 0.0] for _ in range(9)]
        for _, n, i, val in assigns:
            arr[int(n)][int(i)] = float(val)
        checks = re.findall(r"ISCLOSE\(\w\((\d+)\), ([^)]+(?:\))?)\);", sec)
        expected = []
        for k, rhs in checks:
            mm = re.match(r"v\((\d+), (\d+)\)", rhs.strip())
            expected.append([int(k), [int(mm.group(1)), int(mm.group(2))]] if mm else [int(k), float(rhs)])
        vec[name] = {"input": arr, "expected": expected}
    fx["Vector_periodic_Regular_2_2"] = vec

    src = open(os.path.join(basic, "VectorPartitioned.cpp")).read()
    sec = section(src, "basic")
    m = re.search(r"xt::xtensor<bool, 2> is_p = \{", sec)
    lit = _braced(sec, m.end() - 1).replace("true", "True").replace("false", "False")
    fx["VectorPartitioned_is_p_Regular_3_3"] = _cpp_literal(lit)

    src = open(os.path.join(basic, "MatrixDiagonalPartitioned.cpp")).read()
    sec = section(src, "assemble")
    m = re.search(r"xt::xtensor<size_t, 1> iip = \{", sec)
    iip = _cpp_literal(_braced(sec, m.end() - 1))
    m = re.search(r"xt::xtensor<double, 1> a = \{", sec)
    fx["MatrixDiagonalPartitioned_lumped_mass_Regular_2_2"] = {"iip": iip, "data": _cpp_literal(_braced(sec, m.end() - 1))}

    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fixtures.json")
    with open(out, "w") as f:
        json.dump(fx, f, indent=None, separators=(",", ":"))
        f.write("\n")
    print("wrote", out, "keys:", sorted(fx))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
