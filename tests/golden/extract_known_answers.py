"""Extract the known-answer values of the reference's unit tests into
known_answers.json (run once in a container holding the reference; the tests
only read the JSON).

Sources (jdesmarais/dim2D_finite_volume, src/test_files/):
  test_lernaeanhydra/test_dim2d_fluxes.f:303-365        26 DIM flux values at (3,3)
  test_augeanstables/test_mattsson_operators.f:135-152   interior Mattsson stencil values
  test_augeanstables/test_mattsson_operators.f:1580-1637 the 7x5 density field of that test
  test_augeanstables/test_cg_operators.f:113-187          Cockburn-Gau stencil values and field
  test_bc_operators/test_wall_xy_equilibrium.f           wall equilibrium values
  test_bc_operators/test_ridders_method.f                Ridders roots

    python tests/golden/extract_known_answers.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
T = os.path.join(REF, "src", "test_files")
HERE = os.path.dirname(os.path.abspath(__file__))


def read(rel):
    with open(os.path.join(T, rel)) as f:
        return f.read().splitlines()


def num(s):
    return float(s.replace("d", "e").replace("D", "e"))


def flux_test():
    lines = read("test_lernaeanhydra/test_dim2d_fluxes.f")
    text = "\n".join(lines)
    # nodes = reshape((/ ... /), (/5,5,ne/))
    block = text[text.index("nodes = reshape((/"):text.index("(/5,5,ne/))")]
    vals = [num(v) for v in re.findall(r"-?\d+\.\d*d0", block)]
    assert len(vals) == 100, len(vals)
    expected = {}
    for ln in lines:
        m = re.match(r"\s*test_data\((\d+)\)\s*=\s*([-\d.]+d0)", ln)
        if m:
            expected[int(m.group(1))] = num(m.group(2))
    names = []
    for ln in lines:
        m = re.search(r"'(flux_[xy]_[a-z_]+)'\)", ln)
        if m:
            names.append(m.group(1))
    return {"source": "test_lernaeanhydra/test_dim2d_fluxes.f:303-365",
            "nodes_5x5x4_fortran_order": vals, "dx": 0.5, "dy": 0.6, "i": 3, "j": 3,
            "expected": [expected[k] for k in range(1, 27)], "labels": names}


def stencil_test():
    lines = read("test_augeanstables/test_mattsson_operators.f")
    text = "\n".join(lines)
    start = text.index("test interior")
    seg = text[start:text.index("call test_operator", start)]
    exp = [num(m) for m in re.findall(r"test_data\(\d+\)\s*=\s*([-\d.]+d0)", seg)]
    assert len(exp) == 16
    init = text[text.index("subroutine initialize_nodes_x"):text.index("end subroutine initialize_nodes_x")]
    field = {}
    for m in re.finditer(r"nodes\((\d),(\d),1\)=([-\d.]+)", init):
        field[(int(m.group(1)), int(m.group(2)))] = float(m.group(3))
    rows = [[field[(i, j)] for i in range(1, 8)] for j in range(1, 6)]
    return {"source": "test_augeanstables/test_mattsson_operators.f:135-152,1580-1637",
            "note": "default-kind (single precision) literals: inputs and dx, dy are float32 values",
            "rho_rows_j1_to_j5": rows, "dx": 0.5, "dy": 0.6, "i": 3, "j": 3,
            "expected": exp,
            "labels": ["f", "dfdx", "dfdy", "d2fdx2", "d2fdy2", "d2fdxdy", "gradient_x", "g",
                       "dgdx", "dgdy", "d2gdx2", "d2gdy2", "d2gdxdy", "gradient_y", "dfdx_nl",
                       "dgdy_nl"]}


def cg_stencil_test():
    lines = read("test_augeanstables/test_cg_operators.f")
    text = "\n".join(lines)
    start = text.index("test interior")
    seg = text[start:text.index("call test_operator", start)]
    exp = [num(m) for m in re.findall(r"test_data\(\d+\)\s*=\s*([-\d.]+d0)", seg)]
    assert len(exp) == 16
    call = re.search(r"call test_operator\(\s*\$\s*sd_interior,[^)]*?(\d),(\d),\s*(\d),(\d),", text[start:],
                     re.S)
    i_f, j_f, i_g, j_g = (int(call.group(k)) for k in range(1, 5))
    assert (i_f, j_f) == (i_g, j_g)
    init = text[:start]
    field = {}
    for m in re.finditer(r"nodes\((\d),(\d),1\)=([-\d.]+)", init):
        field[(int(m.group(1)), int(m.group(2)))] = float(m.group(3))
    rows = [[field[(i, j)] for i in range(1, 8)] for j in range(1, 6)]
    dx = num(re.search(r"dx\s*=\s*([-\d.]+)", init).group(1))
    dy = num(re.search(r"dy\s*=\s*([-\d.]+)", init).group(1))
    return {"source": "test_augeanstables/test_cg_operators.f:113-187",
            "note": "default-kind (single precision) literals: inputs and dx, dy are float32 values",
            "rho_rows_j1_to_j5": rows, "dx": dx, "dy": dy, "i": i_f, "j": j_f,
            "expected": exp,
            "labels": ["f", "dfdx", "dfdy", "d2fdx2", "d2fdy2", "d2fdxdy", "gradient_x", "g",
                       "dgdx", "dgdy", "d2gdx2", "d2gdy2", "d2gdxdy", "gradient_y", "dfdx_nl",
                       "dgdy_nl"]}


def wall_test():
    text = "\n".join(read("test_bc_operators/test_wall_xy_equilibrium.f"))

    def after(tag):
        i = text.index(tag) + len(tag)
        return num(re.search(r"([-\d.]+d0)", text[i:i + 600]).group(1))
    return {
        "source": "test_bc_operators/test_wall_xy_equilibrium.f",
        "temperature": {"args": [1.5, 50.0, -0.1, 0.2, 6.3],
                        "expected": after("temperature(1.5d0,50.0d0,-0.1d0,0.2d0,6.3d0)")},
        "dwallInternalEnergy_dmd": {"args": [1.5, 0.95, 45.0],
                                    "expected": after("dwallInternalEnergy_dmd(1.5d0,0.95d0,45.0d0)")},
        "wall_x_equilibrium_root_fct": {
            "md": 1.5, "md_x": [0.2, 0.5], "md_y": [1.2, 0.9], "velocity": [0.02, 0.03],
            "Ed": 2.26, "delta": [0.5, 0.8], "angle": 45.0, "heat_flux": 0.005,
            "expected": num(re.search(r"(-0\.2576\d+d0)", text).group(1)),
            "note": "the test's expected value holds for temperature_average = T + "
                    "0.5*delta_x*Pr*q (the test predates the Re*3/(8cv) factor of "
                    "wall_xy_equilibrium_module.f:2295-2336); the pin rescales q accordingly"},
        "brackets": {"fct2": {"T": 0.95, "md_guess": 0.0, "expected": "[md_vap(0.95), md_liq(0.95)]"},
                     "fct4": {"T": 0.95, "md_guess": 1.5, "expected": [1.5, 2.01]}},
    }


def ridders_test():
    return {"source": "test_bc_operators/test_ridders_method.f",
            "cases": [{"fct": "1+x", "x1": -2.0, "x2": 0.0, "root": -1.0},
                      {"fct": "1+x", "x1": -2.0, "x2": 1.0, "root": -1.0},
                      {"fct": "(x-1)(x-2)", "x1": -3.0, "x2": 1.5, "root": 1.0},
                      {"fct": "tanh(x)", "x1": -3.0, "x2": 2.0, "root": 0.0}],
            "tolerance": 1e-10}


if __name__ == "__main__":
    out = {"fluxes": flux_test(), "mattsson": stencil_test(), "cockburn_gau": cg_stencil_test(),
           "wall": wall_test(), "ridders": ridders_test()}
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "known_answers.json"))
